// DfPlan: setup (integrals, metric factor, fitted tensor), Fock builds and the
// device-resident RHF SCF.
//
// Setup follows the reference's prepared CUDA plan
// (src/scf/fock_prepared.cpp:216-330, src/scf/cuda/df_plan_setup.cpp:456-830)
// with an MI355X-specific factorisation: the (P|Q) metric is Cholesky-factored
// in FP64 (rocSOLVER) and B = A L^-T is formed in place by a triangular solve.
// At full rank this gives the same J and K as the reference's M^-1/2 (B B^T =
// A M^-1 A^T); a metric whose smallest Cholesky pivot falls below
// 10 x threshold x max(M_ii) takes the reference's eigen-cutoff route
// (src/integrals/density_fitting_metric.cpp:101-152) instead.
//
// The SCF mirrors run_rhf_host_plan (src/scf/solver/mean_field_driver.cpp:
// 233-345) but keeps every matrix on the device: per iteration only the
// energy and the density RMS (two doubles) cross to the host.
#include "plan.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <numeric>
#include <stdexcept>

#include "scf.hpp"

namespace vqc {
namespace {

constexpr std::size_t kTrsmColumns = 1 << 16;

struct Event {
  hipEvent_t e = nullptr;
  Event() { VQC_HIP(hipEventCreate(&e)); }
  ~Event() {
    if (e) (void)hipEventDestroy(e);
  }
  Event(const Event&) = delete;
  Event& operator=(const Event&) = delete;
  Event(Event&& o) noexcept : e(std::exchange(o.e, nullptr)) {}
};

// Opt-in phase trace: with VIBEQC_TRACE set, one JSON line per phase on stderr
// (the role of the reference's GENERATIVEQC_DF_TRACE, cuda_component_trace.cpp).
class Trace {
 public:
  explicit Trace(const char* scope) : scope_(scope), on_(std::getenv("VIBEQC_TRACE") != nullptr) {
    last_ = start_ = std::chrono::steady_clock::now();
  }
  void mark(const char* phase) {
    if (!on_) return;
    const auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "{\"trace\": \"%s\", \"phase\": \"%s\", \"ms\": %.3f}\n", scope_, phase,
                 std::chrono::duration<double, std::milli>(now - last_).count());
    last_ = now;
  }
  void finish(int n, int naux, std::size_t npair) {
    if (!on_) return;
    std::fprintf(stderr, "{\"trace\": \"%s\", \"phase\": \"total\", \"ms\": %.3f, \"nbf\": %d, \"naux\": %d, \"npair\": %zu}\n",
                 scope_, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start_).count(), n,
                 naux, npair);
  }

 private:
  const char* scope_;
  bool on_;
  std::chrono::steady_clock::time_point start_, last_;
};

// Zero the strict upper triangle of a column-major n x n matrix.
__global__ void zero_strict_upper_kernel(double* a, int n) {
  const std::size_t total = static_cast<std::size_t>(n) * n;
  for (std::size_t k = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x; k < total;
       k += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    const std::size_t row = k % n, col = k / n;
    if (row < col) a[k] = 0.0;
  }
}

// n x n doubles held besides the fitted tensor by a plan and its SCF: S, H,
// J, K, D, the Fock matrix, residual temporaries and an eight-deep DIIS
// history of Fock and error matrices.
constexpr std::size_t kSquareMatrices = 32;

void check_basis_limits(const System& sys, int lmax, const char* what) {
  for (const auto& s : sys.shells)
    if (static_cast<int>(s.l) > lmax)
      throw NotImplemented(std::string(what) + " shells above l=" + std::to_string(lmax) +
                           " are not supported by the device kernels");
}

}  // namespace

namespace {
// gfx architecture name of a device (cached: plans of a batch are created by
// the hundred).
std::string device_arch(int device) {
  static std::mutex mutex;
  static std::map<int, std::string> names;
  std::lock_guard<std::mutex> lock(mutex);
  auto it = names.find(device);
  if (it != names.end()) return it->second;
  hipDeviceProp_t prop{};
  VQC_HIP(hipGetDeviceProperties(&prop, device));
  std::string name(prop.gcnArchName);
  name = name.substr(0, name.find(':'));
  if (name.empty()) name = "unknown";
  names[device] = name;
  return name;
}
}  // namespace

DfPlan::DfPlan(int device, const System& orbital, const System& auxiliary, const PlanOptions& options,
               hipStream_t stream, rocblas_handle blas)
    : device_(device), orb_(orbital), opts_(options) {
  VQC_HIP(hipSetDevice(device_));
  arch_ = device_arch(device_);
  if (stream && blas) {
    stream_ = stream;
    blas_ = blas;
    owns_stream_ = false;
  } else {
    VQC_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    VQC_BLAS(rocblas_create_handle(&blas_));
    VQC_BLAS(rocblas_set_stream(blas_, stream_));
    // Deterministic library kernels: no atomics-based reductions (rocSOLVER's
    // eigensolvers call rocBLAS through this handle too).
    VQC_BLAS(rocblas_set_atomics_mode(blas_, rocblas_atomics_not_allowed));
  }
  try {
    setup(auxiliary);
  } catch (...) {
    if (owns_stream_) {
      rocblas_destroy_handle(blas_);
      (void)hipStreamDestroy(stream_);
    }
    throw;
  }
}

DfPlan::~DfPlan() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (!owns_stream_) return;
  if (blas_) rocblas_destroy_handle(blas_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void DfPlan::sync() { VQC_HIP(hipStreamSynchronize(stream_)); }

std::size_t DfPlan::device_bytes() const {
  return boys_.bytes() + S_.bytes() + H_.bytes() + B64_.bytes() + B32_.bytes() + dpacked_.bytes() +
         rho_.bytes() + rho_partial_.bytes() + J_.bytes() + K_.bytes() + syrk_partial_.bytes() + D_.bytes() +
         c64_.bytes() + c32_.bytes() + ct_hi_.bytes() + ct_lo_.bytes() + U_.bytes() + L_.bytes() + G_.bytes() +
         Ablk_.bytes() + Bblk_.bytes() + Jtmp_.bytes() + Ktmp_.bytes() + Ublk_.bytes() + Kref_.bytes() + Cref_.bytes() +
         facA_.bytes() + facB_.bytes();
}

std::size_t DfPlan::resident_bytes() const {
  const std::size_t nn = static_cast<std::size_t>(n_) * n_, mm = static_cast<std::size_t>(naux_) * naux_;
  const std::size_t doubles = npair_ * ldq_ + 2 * mm + static_cast<std::size_t>(n_) * std::max(nocc_, 1) * ldq_ +
                              std::max(coulomb_partial_size(npair_, ldq_), static_cast<std::size_t>(n_) * ldq_) +
                              exchange_partial_size(n_) + kSquareMatrices * nn;
  return doubles * sizeof(double);
}

void DfPlan::require_resident(const char* what) const {
  if (streamed_)
    throw NotImplemented(std::string(what) +
                         " needs the resident fitted tensor; this plan regenerates it within its device budget");
}

void DfPlan::setup(const System& aux) {
  check_basis_limits(orb_, kMaxOrbitalL, "orbital");
  check_basis_limits(aux, kMaxAuxiliaryL, "auxiliary");
  n_ = static_cast<int>(orb_.nbf());
  naux_ = static_cast<int>(aux.nbf());
  if (n_ <= 0 || naux_ <= 0) throw std::invalid_argument("empty orbital or auxiliary basis");
  ldq_ = static_cast<int>(round_up(naux_, 128));
  npair_ = pair_count(n_);
  enuc_ = orb_.nuclear_repulsion();
  nocc_ = orb_.electrons / 2;
  flag_.resize(1);

  Trace trace("plan_setup");
  boys_.upload(boys_table_host(), stream_);
  aux_ = aux;
  DeviceBasis& ob = ob_;
  upload_basis(orb_, ob, false, stream_);
  AuxiliaryLayout& al = al_;
  build_auxiliary_layout(aux, al, stream_);
  trace.mark("basis_upload");

  const std::size_t nn = static_cast<std::size_t>(n_) * n_;
  S_.resize(nn);
  H_.resize(nn);
  {
    DeviceArray<double> T(nn), V(nn);
    compute_one_electron(orb_, ob, boys_.data(), S_.data(), T.data(), V.data(), stream_);
    assemble_fock(T.data(), V.data(), V.data(), 1.0, 0.0, nn, H_.data(), stream_);
    sync();
  }
  trace.mark("one_electron");
  // VIBEQC_TRACE=checksum: fixed-order sums of squares of the setup's
  // intermediate and final tensors, keyed by the nuclear repulsion (a
  // debugging aid that shows whether two runs built identical tensors).
  const char* trace_env = std::getenv("VIBEQC_TRACE");
  const bool checksums = trace_env && std::strcmp(trace_env, "checksum") == 0;
  auto sum_squares = [&](const double* x, std::size_t count) {
    DeviceArray<double> sums(2);
    dot_pair(x, x, x, count, sums.data(), stream_);
    double h[2];
    sums.download(h, 2, stream_);
    sync();
    return h[0];
  };
  // Device copies of the inputs (uploaded above): Boys table and both bases.
  const double sum_in = checksums ? sum_squares(boys_.data(), boys_.size()) + sum_squares(al.basis.exps.data(), al.basis.exps.size()) +
                                        sum_squares(al.basis.coefs.data(), al.basis.coefs.size()) +
                                        sum_squares(al.hermite.data(), al.hermite.size()) +
                                        sum_squares(ob.exps.data(), ob.exps.size()) + sum_squares(ob.coefs.data(), ob.coefs.size())
                                  : 0.0;
  DeviceArray<double> M(static_cast<std::size_t>(naux_) * naux_);
  compute_metric(al, boys_.data(), M.data(), stream_);
  sync();
  const double sum_m = checksums ? sum_squares(M.data(), M.size()) : 0.0;
  const double sum_in2 = checksums ? sum_squares(boys_.data(), boys_.size()) + sum_squares(al.basis.exps.data(), al.basis.exps.size()) +
                                         sum_squares(al.basis.coefs.data(), al.basis.coefs.size()) +
                                         sum_squares(al.hermite.data(), al.hermite.size()) +
                                         sum_squares(ob.exps.data(), ob.exps.size()) + sum_squares(ob.coefs.data(), ob.coefs.size())
                                   : 0.0;
  const double sum_m2 = checksums ? sum_squares(M.data(), M.size()) : 0.0;
  trace.mark("metric");
  streamed_ = opts_.device_budget > 0 && resident_bytes() > opts_.device_budget;
  if (streamed_) {
    // Keep only the metric factor; B is regenerated per Fock build.
    if (2 * M.bytes() + kSquareMatrices * nn * sizeof(double) > opts_.device_budget)
      throw OutOfMemory("DF memory budget cannot hold the metric and one contraction tile");
    factor_metric(M, nullptr);
    M.release();
    plan_streaming();
    trace.mark(cholesky_ ? "cholesky_inverse" : "eigen_cutoff");
    trace.finish(n_, naux_, npair_);
    return;
  }
  B64_.resize(npair_ * ldq_);
  compute_three_center(orb_, ob, al, boys_.data(), B64_.data(), ldq_, opts_.screening * 1e-10, stream_);
  sync();
  const double sum_a = checksums ? sum_squares(B64_.data(), B64_.size()) : 0.0;
  trace.mark("three_center");
  factor_metric(M, B64_.data());
  trace.mark(cholesky_ ? "cholesky_trsm" : "eigen_cutoff_gemm");
  VQC_HIP(hipMemsetAsync(flag_.data(), 0, sizeof(int), stream_));
  check_finite(B64_.data(), B64_.size(), flag_.data(), stream_);
  int bad = 0;
  flag_.download(&bad, 1, stream_);
  sync();
  trace.mark("finite_check");
  trace.finish(n_, naux_, npair_);
  if (bad) throw NumericalFailure("non-finite fitted three-centre tensor");
  if (checksums)
    std::fprintf(stderr,
                 "{\"trace\": \"setup_checksum\", \"enuc\": %.17g, \"B\": %.17g, \"S\": %.17g, \"H\": %.17g, "
                 "\"L\": %.17g, \"M\": %.17g, \"A\": %.17g, \"cholesky\": %d, \"IN\": %.17g, \"IN2\": %.17g, \"M2\": %.17g}\n",
                 enuc_, sum_squares(B64_.data(), B64_.size()), sum_squares(S_.data(), nn), sum_squares(H_.data(), nn),
                 L_.size() ? sum_squares(L_.data(), L_.size()) : 0.0, sum_m, sum_a, cholesky_ ? 1 : 0, sum_in, sum_in2,
                 sum_m2);
}

void DfPlan::factor_metric(DeviceArray<double>& M, double* A) {
  const int na = naux_;
  const std::size_t mm = static_cast<std::size_t>(na) * na;
  std::vector<double> mdiag(na), ldiag(na);
  DeviceArray<double> diag(na);
  extract_diagonal(M.data(), na, diag.data(), stream_);
  diag.download(mdiag.data(), na, stream_);
  DeviceArray<double> L(mm);
  copy_device(L.data(), M.data(), mm, stream_);
  DeviceArray<rocblas_int> info(1);
  VQC_BLAS(rocsolver_dpotrf(blas_, rocblas_fill_lower, na, L.data(), na, info.data()));
  rocblas_int hinfo = 0;
  info.download(&hinfo, 1, stream_);
  extract_diagonal(L.data(), na, diag.data(), stream_);
  diag.download(ldiag.data(), na, stream_);
  sync();
  const double maxdiag = *std::max_element(mdiag.begin(), mdiag.end());
  if (!(maxdiag > 0.0) || !std::isfinite(maxdiag)) throw NumericalFailure("Coulomb metric has no positive diagonal");
  double minpiv = std::numeric_limits<double>::infinity(), maxpiv = 0.0;
  for (double d : ldiag) {
    minpiv = std::min(minpiv, d * d);
    maxpiv = std::max(maxpiv, d * d);
  }
  const bool use_cholesky = hinfo == 0 && std::isfinite(minpiv) && minpiv > 10.0 * opts_.metric_threshold * maxdiag;
  const double one = 1.0, zero = 0.0;
  if (use_cholesky) {
    cholesky_ = true;
    rank_ = na;
    condition_ = maxpiv / minpiv;
    if (!A) {
      // Streamed: G = L^-1 (lower), so B^T = G A^T block by block.
      hipLaunchKernelGGL(zero_strict_upper_kernel, dim3(1024), dim3(256), 0, stream_, L.data(), na);
      VQC_LAUNCH("zero_strict_upper_kernel");
      VQC_BLAS(rocsolver_dtrtri(blas_, rocblas_fill_lower, rocblas_diagonal_non_unit, na, L.data(), na, info.data()));
      info.download(&hinfo, 1, stream_);
      sync();
      if (hinfo != 0) throw NumericalFailure("metric Cholesky factor is singular");
      G_ = std::move(L);
      g_lower_ = true;
      return;
    }
    for (std::size_t c0 = 0; c0 < npair_; c0 += kTrsmColumns) {
      const int cols = static_cast<int>(std::min(kTrsmColumns, npair_ - c0));
      VQC_BLAS(rocblas_dtrsm(blas_, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                             rocblas_diagonal_non_unit, na, cols, &one, L.data(), na, A + c0 * ldq_, ldq_));
    }
    sync();
    L_ = std::move(L);  // kept for the gradient's c = L^-T rho and Y = L^-T T
    return;
  }
  // Eigen-cutoff route: X = V diag(lambda^-1/2 | lambda > thr lambda_max) V^T, B = A X.
  cholesky_ = false;
  DeviceArray<double> V(mm), w(na), e(na);
  copy_device(V.data(), M.data(), mm, stream_);
  VQC_BLAS(rocsolver_dsyevd(blas_, rocblas_evect_original, rocblas_fill_lower, na, V.data(), na, w.data(), e.data(),
                            info.data()));
  std::vector<double> hw(na);
  w.download(hw.data(), na, stream_);
  info.download(&hinfo, 1, stream_);
  sync();
  if (hinfo != 0) throw NumericalFailure("metric eigensolver did not converge");
  const double largest = hw.back();
  if (!(largest > 0.0)) throw NumericalFailure("Coulomb metric has no positive eigenspace");
  const double cutoff = opts_.metric_threshold * largest;
  std::vector<double> scale(na, 0.0);
  rank_ = 0;
  double smallest = largest;
  for (int k = 0; k < na; ++k)
    if (hw[k] > cutoff) {
      scale[k] = 1.0 / std::sqrt(hw[k]);
      ++rank_;
      smallest = std::min(smallest, hw[k]);
    }
  if (rank_ == 0) throw NumericalFailure("metric threshold removed every auxiliary direction");
  condition_ = largest / smallest;
  // Scale columns: Vs[:, k] = V[:, k] * scale[k]  (column-major, scale applied as w^-1/2 of 1/scale^2).
  std::vector<double> inv(na);
  for (int k = 0; k < na; ++k) inv[k] = scale[k] > 0.0 ? 1.0 / (scale[k] * scale[k]) : std::numeric_limits<double>::infinity();
  DeviceArray<double> winv(na), Vs(mm), X(mm);
  winv.upload(inv, stream_);
  scale_columns_inv_sqrt(V.data(), winv.data(), na, Vs.data(), stream_);
  VQC_BLAS(rocblas_dgemm(blas_, rocblas_operation_none, rocblas_operation_transpose, na, na, na, &one, Vs.data(), na,
                         V.data(), na, &zero, X.data(), na));
  if (!A) {
    sync();
    G_ = std::move(X);
    g_lower_ = false;
    return;
  }
  DeviceArray<double> tmp(kTrsmColumns * ldq_);
  VQC_HIP(hipMemsetAsync(tmp.data(), 0, tmp.bytes(), stream_));
  for (std::size_t c0 = 0; c0 < npair_; c0 += kTrsmColumns) {
    const int cols = static_cast<int>(std::min(kTrsmColumns, npair_ - c0));
    VQC_BLAS(rocblas_dgemm(blas_, rocblas_operation_none, rocblas_operation_none, na, cols, na, &one, X.data(), na,
                           A + c0 * ldq_, ldq_, &zero, tmp.data(), ldq_));
    copy_device(A + c0 * ldq_, tmp.data(), static_cast<std::size_t>(cols) * ldq_, stream_);
  }
  sync();
}

void DfPlan::set_precision(Precision p) {
  // A streamed plan regenerates B in FP64 for every build and has no reduced copy.
  prec_ = streamed_ ? Precision::FP64 : p;
  ensure_reduced();
}

void DfPlan::ensure_reduced() {
  if (prec_ == Precision::FP64 || B32_.size() || streamed_) return;
  B32_.resize(npair_ * ldq_);
  convert_f64_to_f32(B64_.data(), B32_.data(), B32_.size(), stream_);
  sync();
}

void DfPlan::ensure_factor(int r) {
  const std::size_t nn = static_cast<std::size_t>(n_) * n_;
  if (!J_.size()) {
    J_.resize(nn);
    K_.resize(nn);
    D_.resize(nn);
    dpacked_.resize(npair_);
    rho_.resize(ldq_);
    rho_partial_.resize(std::max(coulomb_partial_size(npair_, ldq_), static_cast<std::size_t>(n_) * ldq_));
    // Streamed plans bound the split-K slices (partial_capacity) to ten n x n matrices.
    syrk_partial_.resize(streamed_ ? std::min(exchange_partial_size(n_), 10 * nn) : exchange_partial_size(n_));
  }
  r = std::max(r, 1);
  const JkShape shape = shape_for(prec_, r);
  const int ni = shape.u_tile > 0 ? shape.u_tile : occupied_tile(r);
  const int ldc = static_cast<int>(round_up(r, 32 * ni));
  if (r <= factor_cap_ && ni == ni_ && ldc == ldc_) return;
  ni_ = ni;
  ldc_ = ldc;
  factor_cap_ = r;
  const int ldct = static_cast<int>(round_up(n_, 32));
  c64_.resize(static_cast<std::size_t>(n_) * ldc_);
  c32_.resize(static_cast<std::size_t>(n_) * ldc_);
  ct_hi_.resize(static_cast<std::size_t>(ldc_) * ldct);
  ct_lo_.resize(static_cast<std::size_t>(ldc_) * ldct);
  U_.release();
  // FP64 U (r ldq doubles per mu), or the two BF16X3 U of an incremental
  // build (2 x 4 (r ldq + 64) bytes per mu).  Streamed plans: Ublk_ per block.
  if (!streamed_) U_.resize(static_cast<std::size_t>(n_) * (static_cast<std::size_t>(r) * ldq_ + 64) * sizeof(double));
  ops_ = OccupiedOperands{};
  ops_.n = n_;
  ops_.ldc = ldc_;
  ops_.ni = ni_;
  ops_.ldct = ldct;
  ops_.c64 = c64_.data();
  ops_.c32 = c32_.data();
  ops_.ct_hi = ct_hi_.data();
  ops_.ct_lo = ct_lo_.data();
}

void DfPlan::build_jk(Precision p, const double* D, int r, int r_pos) {
  if (streamed_) {
    build_jk_streamed(D, r, r_pos, opts_.exchange);
    return;
  }
  if (p != Precision::FP64) ensure_reduced();
  const void* B = p == Precision::FP64 ? static_cast<const void*>(B64_.data()) : static_cast<const void*>(B32_.data());
  // When the U pass runs it also produces the rho partials, so J needs only
  // one more pass over B.
  const bool fused_rho = opts_.exchange && opts_.coulomb && r > 0;
  if (opts_.exchange) {
    ops_.r = r;
    ops_.ustride = u_row_stride(p, std::max(r, 1), ldq_);
    ops_.terms = p == Precision::BF16X3 ? bf16_terms_ : 3;
    if (r > 0) {
      prepare_occupied(p, ops_, stream_);
      occupied_u(p, B, ops_, ldq_, U_.data(), D, fused_rho ? rho_partial_.data() : nullptr, stream_);
    }
    exchange_syrk(p, U_.data(), ops_.ustride, n_, r, r_pos, ldq_, K_.data(), syrk_partial_.data(), stream_, 0,
                  shape_for(p, std::max(r, 1)).syrk_splits, ops_.terms);
    ops_.terms = 3;
  }
  if (opts_.coulomb) {
    if (fused_rho) {
      rho_from_partials(rho_partial_.data(), n_, ldq_, rho_.data(), stream_);
    } else {
      pack_density(D, n_, dpacked_.data(), stream_);
      coulomb_rho(p, B, dpacked_.data(), npair_, ldq_, rho_.data(), rho_partial_.data(), stream_);
    }
    coulomb_matrix(p, B, rho_.data(), npair_, ldq_, n_, J_.data(), stream_);
  }
}

bool DfPlan::incremental_enabled() const {
  if (streamed_ || !opts_.exchange || prec_ != Precision::BF16X3) return false;
  const char* v = std::getenv("VIBEQC_INCREMENTAL_K");
  return !(v && v[0] == '0');
}

namespace {
// A = (C + R) / 2, B = C - R over `count` elements.
__global__ void half_sum_difference_kernel(const double* __restrict__ C, const double* __restrict__ R,
                                           std::size_t count, double* __restrict__ A, double* __restrict__ B) {
  const std::size_t stride = static_cast<std::size_t>(gridDim.x) * blockDim.x;
  for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
    const double c = C[i], r = R[i];
    A[i] = 0.5 * (c + r);
    B[i] = c - r;
  }
}
}  // namespace

void DfPlan::build_jk_incremental(int r) {
  ensure_reduced();
  const std::size_t nn = static_cast<std::size_t>(n_) * n_;
  const std::size_t fc = static_cast<std::size_t>(n_) * ldc_;
  if (Kref_.size() != nn || Cref_.size() != fc) throw std::logic_error("incremental build without an anchor");
  facA_.resize(fc);
  facB_.resize(fc);
  // J of the current density in FP64 (one pass over B for rho, one for J).
  coulomb_only(Precision::FP64, D_.data());
  hipLaunchKernelGGL(half_sum_difference_kernel, dim3(static_cast<unsigned>(std::min<std::size_t>((fc + 255) / 256, 4096))),
                     dim3(256), 0, stream_, c64_.data(), Cref_.data(), fc, facA_.data(), facB_.data());
  VQC_LAUNCH("half_sum_difference_kernel");
  const std::size_t ustride = u_row_stride(Precision::BF16X3, r, ldq_);
  char* ua = U_.data();
  char* ub = ua + static_cast<std::size_t>(n_) * ustride * 2 * sizeof(uint16_t);
  if (static_cast<std::size_t>(ub - ua) * 2 > U_.size()) throw std::logic_error("U buffer too small for two BF16X3 U");
  OccupiedOperands op = ops_;
  op.r = r;
  op.ustride = ustride;
  op.c64 = facA_.data();
  prepare_occupied(Precision::BF16X3, op, stream_);
  occupied_u(Precision::BF16X3, B32_.data(), op, ldq_, ua, nullptr, nullptr, stream_);
  op.c64 = facB_.data();
  prepare_occupied(Precision::BF16X3, op, stream_);
  occupied_u(Precision::BF16X3, B32_.data(), op, ldq_, ub, nullptr, nullptr, stream_);
  exchange_cross_bf16x3(ua, ub, ustride, n_, r, ldq_, Kref_.data(), K_.data(), syrk_partial_.data(), stream_);
}

// Streamed plans (the reference's budgeted DF planner,
// src/scf/density_fitting.cpp:1041-1218, regenerates B in tiles when the full
// tensor does not fit).  Here the fitted tensor is produced one auxiliary block
// B[:, q0:q1] at a time: the three-centre rows are regenerated tile by tile
// (shell ranges) and mapped by one GEMM, B^T[q0:q1, rows] = G[q0:q1, :] A^T.
// J and K are sums over Q, so every block adds its own rho_Q, J and K
// contributions; the block order is fixed, so results are reproducible.
void DfPlan::plan_streaming() {
  const std::size_t budget = opts_.device_budget;
  const std::size_t nn = static_cast<std::size_t>(n_) * n_;
  const std::size_t fixed =
      G_.bytes() + boys_.bytes() +
      (kSquareMatrices * nn + std::min(exchange_partial_size(n_), 10 * nn) + ldq_ + npair_ +
       std::max(coulomb_partial_size(npair_, ldq_), static_cast<std::size_t>(n_) * ldq_)) *
          sizeof(double);
  if (fixed >= budget) throw OutOfMemory("DF memory budget cannot hold the metric and one contraction tile");
  const std::size_t avail = budget - fixed;
  // Row tiles: consecutive shell ranges whose packed rows fit a quarter of what is left.
  const std::size_t row_cap = avail / 4 / (static_cast<std::size_t>(ldq_) * sizeof(double));
  const auto& shells = ob_.host_shells;
  const int nshell = static_cast<int>(shells.size());
  auto first_fn = [&](int sidx) {
    return sidx < nshell ? static_cast<std::size_t>(shells[sidx].fn_off) : static_cast<std::size_t>(n_);
  };
  auto rows = [&](int s0, int s1) { return pair_count(first_fn(s1)) - pair_count(first_fn(s0)); };
  tile_shells_.assign(1, 0);
  std::size_t max_rows = 0;
  for (int s0 = 0; s0 < nshell;) {
    int s1 = s0 + 1;
    while (s1 < nshell && rows(s0, s1 + 1) <= row_cap) ++s1;
    if (rows(s0, s1) > row_cap)
      throw OutOfMemory("DF memory budget cannot hold the three-centre rows of one shell");
    max_rows = std::max(max_rows, rows(s0, s1));
    tile_shells_.push_back(s1);
    s0 = s1;
  }
  const std::size_t tile_bytes = max_rows * ldq_ * sizeof(double);
  // Auxiliary block width for the SCF's occupied rank (a wider factor may
  // narrow a build's blocks, see build_jk_streamed).
  const std::size_t per_column = (npair_ + static_cast<std::size_t>(n_) * (std::max(nocc_, 1) + 1)) * sizeof(double);
  const std::size_t q = (avail - tile_bytes) / per_column / 128 * 128;
  if (q < 128) throw OutOfMemory("DF memory budget cannot hold one auxiliary block of the fitted tensor");
  qblk_ = static_cast<int>(std::min<std::size_t>(q, ldq_));
}

void DfPlan::generate_block(int q0, int q1, int ldqb) {
  const double one = 1.0, zero = 0.0;
  VQC_HIP(hipMemsetAsync(Bblk_.data(), 0, npair_ * ldqb * sizeof(double), stream_));
  const int k = g_lower_ ? q1 : naux_;  // G = L^-1 has no entries right of its diagonal
  const auto& shells = ob_.host_shells;
  const int nshell = static_cast<int>(shells.size());
  for (std::size_t t = 0; t + 1 < tile_shells_.size(); ++t) {
    const int s0 = tile_shells_[t], s1 = tile_shells_[t + 1];
    const std::size_t f0 = shells[s0].fn_off, f1 = s1 < nshell ? shells[s1].fn_off : n_;
    const std::size_t row0 = pair_count(f0), rows = pair_count(f1) - row0;
    compute_three_center(orb_, ob_, al_, boys_.data(), Ablk_.data(), ldq_, opts_.screening * 1e-10, stream_, s0, s1);
    VQC_BLAS(rocblas_dgemm(blas_, rocblas_operation_none, rocblas_operation_none, q1 - q0, static_cast<int>(rows), k,
                           &one, G_.data() + q0, naux_, Ablk_.data(), ldq_, &zero, Bblk_.data() + row0 * ldqb, ldqb));
  }
}

void DfPlan::build_jk_streamed(const double* D, int r, int r_pos, bool exchange) {
  const std::size_t nn = static_cast<std::size_t>(n_) * n_;
  const bool do_k = exchange && r > 0;
  const bool fused_rho = do_k && opts_.coulomb;
  // The U block holds at most the planned factor width; a wider factor (an
  // indefinite or high-rank density) is taken in column chunks of that width,
  // each copied into a zero-padded factor of its own.
  const int rc = std::min(r, std::max(nocc_, 1));
  const int qb = qblk_;
  std::size_t max_rows = 0;
  const auto& shells = ob_.host_shells;
  const int nshell = static_cast<int>(shells.size());
  for (std::size_t t = 0; t + 1 < tile_shells_.size(); ++t) {
    const int s1 = tile_shells_[t + 1];
    const std::size_t f0 = shells[tile_shells_[t]].fn_off, f1 = s1 < nshell ? shells[s1].fn_off : n_;
    max_rows = std::max(max_rows, pair_count(f1) - pair_count(f0));
  }
  if (Ablk_.size() < max_rows * ldq_) Ablk_.resize(max_rows * ldq_);
  if (Bblk_.size() < npair_ * qb) Bblk_.resize(npair_ * qb);
  if (Jtmp_.size() != nn) {
    Jtmp_.resize(nn);
    Ktmp_.resize(nn);
  }
  if (do_k && Ublk_.size() < static_cast<std::size_t>(n_) * rc * qb * sizeof(double)) {
    Ublk_.release();
    Ublk_.resize(static_cast<std::size_t>(n_) * rc * qb * sizeof(double));
  }
  // Column chunks of the factor: [c0, c0 + width), the first `positive` of them +1.
  struct Chunk {
    OccupiedOperands ops;
    int positive;
  };
  std::vector<Chunk> chunks;
  DeviceArray<double> chunk_factor;
  if (do_k) {
    if (r <= rc) {
      ops_.r = r;
      prepare_occupied(Precision::FP64, ops_, stream_);
      chunks.push_back({ops_, r_pos});
    } else {
      const int ni = occupied_tile(rc);
      const int ldcc = static_cast<int>(round_up(rc, 32 * ni));
      const int nchunk = (r + rc - 1) / rc;
      chunk_factor.resize(static_cast<std::size_t>(nchunk) * n_ * ldcc);
      VQC_HIP(hipMemsetAsync(chunk_factor.data(), 0, chunk_factor.bytes(), stream_));
      for (int c = 0; c < nchunk; ++c) {
        const int c0 = c * rc, width = std::min(rc, r - c0);
        double* dst = chunk_factor.data() + static_cast<std::size_t>(c) * n_ * ldcc;
        VQC_HIP(hipMemcpy2DAsync(dst, ldcc * sizeof(double), c64_.data() + c0, ldc_ * sizeof(double),
                                 width * sizeof(double), n_, hipMemcpyDeviceToDevice, stream_));
        OccupiedOperands o = ops_;
        o.r = width;
        o.ni = ni;
        o.ldc = ldcc;
        o.c64 = dst;
        chunks.push_back({o, std::clamp(r_pos - c0, 0, width)});
      }
    }
  }
  VQC_HIP(hipMemsetAsync(J_.data(), 0, nn * sizeof(double), stream_));
  VQC_HIP(hipMemsetAsync(K_.data(), 0, nn * sizeof(double), stream_));
  if (opts_.coulomb && !fused_rho) pack_density(D, n_, dpacked_.data(), stream_);
  const double one = 1.0;
  for (int q0 = 0; q0 < naux_; q0 += qb) {
    const int q1 = std::min(q0 + qb, naux_);
    const int ldqb = static_cast<int>(round_up(q1 - q0, 128));
    generate_block(q0, q1, ldqb);
    for (std::size_t c = 0; c < chunks.size(); ++c) {
      Chunk& ch = chunks[c];
      ch.ops.ustride = u_row_stride(Precision::FP64, ch.ops.r, ldqb);
      occupied_u(Precision::FP64, Bblk_.data(), ch.ops, ldqb, Ublk_.data(), D,
                 fused_rho && c == 0 ? rho_partial_.data() : nullptr, stream_);
      exchange_syrk(Precision::FP64, Ublk_.data(), ch.ops.ustride, n_, ch.ops.r, ch.positive, ldqb, Ktmp_.data(),
                    syrk_partial_.data(), stream_, syrk_partial_.size());
      VQC_BLAS(rocblas_daxpy(blas_, static_cast<int>(nn), &one, Ktmp_.data(), 1, K_.data(), 1));
    }
    if (opts_.coulomb) {
      if (fused_rho)
        rho_from_partials(rho_partial_.data(), n_, ldqb, rho_.data(), stream_);
      else
        coulomb_rho(Precision::FP64, Bblk_.data(), dpacked_.data(), npair_, ldqb, rho_.data(), rho_partial_.data(),
                    stream_);
      coulomb_matrix(Precision::FP64, Bblk_.data(), rho_.data(), npair_, ldqb, n_, Jtmp_.data(), stream_);
      VQC_BLAS(rocblas_daxpy(blas_, static_cast<int>(nn), &one, Jtmp_.data(), 1, J_.data(), 1));
    }
  }
}

void DfPlan::upload_factor(const std::vector<double>& occupied, int nocc) {
  ensure_factor(nocc);
  std::vector<double> c(static_cast<std::size_t>(n_) * ldc_, 0.0);
  const double root2 = std::sqrt(2.0);
  for (int nu = 0; nu < n_; ++nu)
    for (int i = 0; i < nocc; ++i)
      c[static_cast<std::size_t>(nu) * ldc_ + i] = root2 * occupied[static_cast<std::size_t>(nu) * nocc + i];
  c64_.upload(c.data(), c.size(), stream_);
  // D = 2 C C^T = Cs Cs^T with Cs = sqrt(2) C; c64_ is the column-major ldc x n matrix Cs^T.
  const double one = 1.0, zero = 0.0;
  VQC_BLAS(rocblas_dgemm(blas_, rocblas_operation_transpose, rocblas_operation_none, n_, n_, std::max(nocc, 1), &one,
                         c64_.data(), ldc_, c64_.data(), ldc_, &zero, D_.data(), n_));
}

int DfPlan::factorise_density(const double* D) {
  // D = V diag(w) V^T; factor columns sqrt|w_k| v_k, positive eigenvalues first.
  const std::size_t nn = static_cast<std::size_t>(n_) * n_;
  DeviceArray<double> V(nn), w(n_), e(n_);
  DeviceArray<rocblas_int> info(1);
  copy_device(V.data(), D, nn, stream_);
  VQC_BLAS(rocsolver_dsyevd(blas_, rocblas_evect_original, rocblas_fill_lower, n_, V.data(), n_, w.data(), e.data(),
                            info.data()));
  std::vector<double> hw(n_);
  rocblas_int hinfo = 0;
  w.download(hw.data(), n_, stream_);
  info.download(&hinfo, 1, stream_);
  sync();
  if (hinfo != 0) throw NumericalFailure("density eigensolver did not converge");
  double largest = 0.0;
  for (double x : hw) largest = std::max(largest, std::abs(x));
  std::vector<int> pos, neg;
  for (int k = n_ - 1; k >= 0; --k)
    if (hw[k] > 1e-14 * largest) pos.push_back(k);
  for (int k = 0; k < n_; ++k)
    if (hw[k] < -1e-14 * largest) neg.push_back(k);
  std::vector<int> idx = pos;
  idx.insert(idx.end(), neg.begin(), neg.end());
  const int r = static_cast<int>(idx.size());
  ensure_factor(r);
  DeviceArray<int> didx;
  if (r) didx.upload(idx, stream_);
  factor_from_eigen(V.data(), w.data(), r ? didx.data() : nullptr, n_, r, ldc_, c64_.data(), stream_);
  sync();
  return static_cast<int>(pos.size()) | (r << 16);
}

std::vector<double> symmetrised_density(const std::vector<double>& d, int n) {
  double scale = 0.0, asym = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const double a = d[static_cast<std::size_t>(i) * n + j], b = d[static_cast<std::size_t>(j) * n + i];
      if (!std::isfinite(a)) throw std::invalid_argument("density must be finite");
      scale = std::max(scale, std::abs(a));
      asym = std::max(asym, std::abs(a - b));
    }
  if (asym > 1e-10 * std::max(1.0, scale))
    throw std::invalid_argument("density-fitted exchange requires a symmetric density");
  std::vector<double> out(d.size());
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      out[static_cast<std::size_t>(i) * n + j] =
          0.5 * (d[static_cast<std::size_t>(i) * n + j] + d[static_cast<std::size_t>(j) * n + i]);
  return out;
}

bool split_density(const std::vector<double>& d, int n, std::vector<double>& sym, std::vector<double>& anti) {
  double scale = 0.0, asym = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const double a = d[static_cast<std::size_t>(i) * n + j], b = d[static_cast<std::size_t>(j) * n + i];
      if (!std::isfinite(a)) throw std::invalid_argument("density must be finite");
      scale = std::max(scale, std::abs(a));
      asym = std::max(asym, std::abs(a - b));
    }
  const bool nonsymmetric = asym > 1e-14 * scale;
  sym.resize(d.size());
  if (nonsymmetric) anti.resize(d.size());
  else anti.clear();
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const double a = d[static_cast<std::size_t>(i) * n + j], b = d[static_cast<std::size_t>(j) * n + i];
      sym[static_cast<std::size_t>(i) * n + j] = 0.5 * (a + b);
      if (nonsymmetric) anti[static_cast<std::size_t>(i) * n + j] = 0.5 * (a - b);
    }
  return nonsymmetric;
}

void DfPlan::add_general_exchange(const double* A, double* K) {
  require_resident("a nonsymmetric density");
  const int n = n_;
  const std::size_t nn = static_cast<std::size_t>(n) * n;
  const double one = 1.0, zero = 0.0;
  // The row-major array of A read column-major is A^T: A^T A = (A_cm)(A_cm)^T.
  DeviceArray<double> S(nn), V(nn), w(n), e(n), AV(nn);
  DeviceArray<rocblas_int> info(1);
  VQC_BLAS(rocblas_dgemm(blas_, rocblas_operation_none, rocblas_operation_transpose, n, n, n, &one, A, n, A, n, &zero,
                         S.data(), n));
  copy_device(V.data(), S.data(), nn, stream_);
  VQC_BLAS(rocsolver_dsyevd(blas_, rocblas_evect_original, rocblas_fill_lower, n, V.data(), n, w.data(), e.data(),
                            info.data()));
  std::vector<double> hw(n);
  rocblas_int hinfo = 0;
  w.download(hw.data(), n, stream_);
  info.download(&hinfo, 1, stream_);
  sync();
  if (hinfo != 0) throw NumericalFailure("density eigensolver did not converge");
  const double largest = std::max(hw.back(), 0.0);
  std::vector<int> idx;
  std::vector<double> sigma(n, 0.0), inv_sigma(n, 0.0);
  for (int k = n - 1; k >= 0; --k)
    if (hw[k] > 1e-28 * largest && hw[k] > 0.0) {
      idx.push_back(k);
      sigma[k] = std::sqrt(hw[k]);
      inv_sigma[k] = 1.0 / sigma[k];
    }
  const int r = static_cast<int>(idx.size());
  if (r == 0) return;
  // A V = (A_cm)^T V: columns A v_k.
  VQC_BLAS(rocblas_dgemm(blas_, rocblas_operation_transpose, rocblas_operation_none, n, n, n, &one, A, n, V.data(), n,
                         &zero, AV.data(), n));
  ensure_factor(r);
  const int ni = occupied_tile(r);
  const int ldc = static_cast<int>(round_up(r, 32 * ni));
  DeviceArray<double> cp(static_cast<std::size_t>(n) * ldc), cq(static_cast<std::size_t>(n) * ldc), ws(n), wi(n);
  DeviceArray<int> didx;
  didx.upload(idx, stream_);
  ws.upload(sigma, stream_);
  wi.upload(inv_sigma, stream_);
  // Columns: P_k = A v_k / sqrt(s_k), Q_k = v_k sqrt(s_k).
  factor_from_eigen(AV.data(), wi.data(), didx.data(), n, r, ldc, cp.data(), stream_);
  factor_from_eigen(V.data(), ws.data(), didx.data(), n, r, ldc, cq.data(), stream_);
  OccupiedOperands op;
  op.n = n;
  op.r = r;
  op.ldc = ldc;
  op.ni = ni;
  op.ldct = static_cast<int>(round_up(n, 32));
  op.ustride = u_row_stride(Precision::FP64, r, ldq_);
  op.c64 = cp.data();
  OccupiedOperands oq = op;
  oq.c64 = cq.data();
  DeviceArray<double> Uq(static_cast<std::size_t>(n) * op.ustride), M(nn);
  occupied_u(Precision::FP64, B64_.data(), op, ldq_, U_.data(), nullptr, nullptr, stream_);
  occupied_u(Precision::FP64, B64_.data(), oq, ldq_, Uq.data(), nullptr, nullptr, stream_);
  // Row-major M[mu][lam] = sum_k U_P[mu][k] U_Q[lam][k]: column-major M^T = U_Q^T U_P.
  const int kd = static_cast<int>(static_cast<std::size_t>(r) * ldq_);
  VQC_BLAS(rocblas_dgemm(blas_, rocblas_operation_transpose, rocblas_operation_none, n, n, kd, &one, Uq.data(),
                         static_cast<int>(op.ustride), static_cast<const double*>(static_cast<const void*>(U_.data())),
                         static_cast<int>(op.ustride), &zero, M.data(), n));
  VQC_BLAS(rocblas_daxpy(blas_, static_cast<int>(nn), &one, M.data(), 1, K, 1));
  sync();
}

JkResult DfPlan::evaluate(const std::vector<double>& density, std::vector<double>* gradient) {
  converged_state_ = false;
  VQC_HIP(hipSetDevice(device_));
  const std::size_t nn = static_cast<std::size_t>(n_) * n_;
  if (density.size() != nn) throw std::invalid_argument("density shape mismatch");
  std::vector<double> sym, anti;
  const bool nonsymmetric = split_density(density, n_, sym, anti);
  if (nonsymmetric && gradient)
    throw std::invalid_argument("the fixed-density derivative requires a symmetric density");
  ensure_factor(1);
  D_.upload(sym.data(), nn, stream_);
  int packed = opts_.exchange ? factorise_density(D_.data()) : 0;
  const int r_pos = packed & 0xffff, r = packed >> 16;
  build_jk(gradient ? Precision::FP64 : prec_, D_.data(), r, r_pos);
  if (nonsymmetric && opts_.exchange) {
    // J depends on the symmetric part only; K gains sum_Q B_Q D_a B_Q.
    DeviceArray<double> Da(nn);
    Da.upload(anti.data(), nn, stream_);
    add_general_exchange(Da.data(), K_.data());
  }
  JkResult out;
  out.J.resize(nn);
  out.K.resize(nn);
  J_.download(out.J.data(), nn, stream_);
  K_.download(out.K.data(), nn, stream_);
  sync();
  if (!opts_.coulomb) std::fill(out.J.begin(), out.J.end(), 0.0);
  if (!opts_.exchange) std::fill(out.K.begin(), out.K.end(), 0.0);
  if (gradient) {
    require_resident("the fixed-density derivative");
    const std::size_t ng = orb_.atoms.size() * 3;
    DeviceArray<double> g(ng);
    VQC_HIP(hipMemsetAsync(g.data(), 0, g.bytes(), stream_));
    df_two_electron_gradient(r, r_pos, g.data());
    gradient->assign(ng, 0.0);
    g.download(gradient->data(), ng, stream_);
    sync();
  }
  return out;
}

JkResult DfPlan::evaluate_occupied(const std::vector<double>& occupied, int nocc) {
  converged_state_ = false;
  VQC_HIP(hipSetDevice(device_));
  if (nocc < 0 || occupied.size() != static_cast<std::size_t>(n_) * nocc)
    throw std::invalid_argument("occupied factor shape mismatch");
  for (double x : occupied)
    if (!std::isfinite(x)) throw std::invalid_argument("occupied factor must be finite");
  ensure_factor(nocc);
  upload_factor(occupied, nocc);
  build_jk(prec_, D_.data(), nocc, nocc);
  const std::size_t nn = static_cast<std::size_t>(n_) * n_;
  JkResult out;
  out.J.resize(nn);
  out.K.resize(nn);
  J_.download(out.J.data(), nn, stream_);
  K_.download(out.K.data(), nn, stream_);
  sync();
  return out;
}

void DfPlan::evaluate_occupied_into(const double* C, int nocc, double* J, double* K, void* scratch,
                                    std::size_t bytes) {
  VQC_HIP(hipSetDevice(device_));
  converged_state_ = false;
  if (nocc < 1 || nocc > n_) throw std::invalid_argument("occupied column count out of range");
  if (!C || !J || !K || !scratch) throw std::invalid_argument("null device buffer");
  require_resident("the allocation-free occupied build");
  const Precision p = prec_;
  if (p != Precision::FP64 && !B32_.size())
    throw std::logic_error("the reduced-precision copy of B is made by set_precision");
  // Workspace carved from the scratch block, each piece 256-byte aligned.
  std::size_t used = 0;
  auto carve = [&](std::size_t size) {
    used = round_up(used, 256);
    if (size > bytes || used > bytes - size)
      throw std::length_error("workspace exceeds the " + std::to_string(bytes >> 20) + " MiB allocation-free scratch");
    void* at = static_cast<char*>(scratch) + used;
    used += size;
    return at;
  };
  OccupiedOperands ops;
  ops.n = n_;
  ops.r = nocc;
  ops.ni = occupied_tile(nocc);
  ops.ldc = static_cast<int>(round_up(nocc, 32 * ops.ni));
  ops.ldct = static_cast<int>(round_up(n_, 32));
  ops.ustride = u_row_stride(p, nocc, ldq_);
  const std::size_t nn = static_cast<std::size_t>(n_) * n_;
  double* c64 = static_cast<double*>(carve(static_cast<std::size_t>(n_) * ops.ldc * sizeof(double)));
  ops.c64 = c64;
  if (p == Precision::FP32) ops.c32 = static_cast<float*>(carve(static_cast<std::size_t>(n_) * ops.ldc * sizeof(float)));
  if (p == Precision::BF16X3) {
    const std::size_t plane = static_cast<std::size_t>(ops.ldc) * ops.ldct * sizeof(uint16_t);
    ops.ct_hi = carve(plane);
    ops.ct_lo = carve(plane);
  }
  double* D = static_cast<double*>(carve(nn * sizeof(double)));
  double* rho_part = static_cast<double*>(carve(static_cast<std::size_t>(n_) * ldq_ * sizeof(double)));
  double* rho = static_cast<double*>(carve(static_cast<std::size_t>(ldq_) * sizeof(double)));
  const std::size_t u_elem = p == Precision::FP64 ? sizeof(double) : sizeof(float);  // BF16X3: two 2-byte planes
  void* U = carve(static_cast<std::size_t>(n_) * ops.ustride * u_elem);
  // The split-K partials take what is left (at least the two slices of one split).
  used = round_up(used, 256);
  if (used > bytes || (bytes - used) / sizeof(double) < 2 * nn)
    throw std::length_error("workspace exceeds the " + std::to_string(bytes >> 20) + " MiB allocation-free scratch");
  double* partial = static_cast<double*>(carve((bytes - used) / sizeof(double) * sizeof(double)));
  const std::size_t partial_capacity = std::min((bytes - (reinterpret_cast<char*>(partial) - static_cast<char*>(scratch))) /
                                                    sizeof(double),
                                                exchange_partial_size(n_));
  const void* B = p == Precision::FP64 ? static_cast<const void*>(B64_.data()) : static_cast<const void*>(B32_.data());
  occupied_factor_and_density(C, n_, nocc, ops.ldc, c64, D, stream_);
  prepare_occupied(p, ops, stream_);
  occupied_u(p, B, ops, ldq_, U, D, rho_part, stream_);
  exchange_syrk(p, U, ops.ustride, n_, nocc, nocc, ldq_, K, partial, stream_, partial_capacity);
  rho_from_partials(rho_part, n_, ldq_, rho, stream_);
  coulomb_matrix(p, B, rho, npair_, ldq_, n_, J, stream_);
  sync();
}

std::array<double, 5> DfPlan::benchmark_occupied(const std::vector<double>& occupied, int nocc, unsigned repeat) {
  converged_state_ = false;
  VQC_HIP(hipSetDevice(device_));
  if (repeat == 0) throw std::invalid_argument("repeat must be positive");
  require_resident("the kernel benchmark");
  if (nocc <= 0 || occupied.size() != static_cast<std::size_t>(n_) * nocc)
    throw std::invalid_argument("occupied factor shape mismatch");
  ensure_factor(nocc);
  upload_factor(occupied, nocc);
  const Precision p = prec_;
  if (p != Precision::FP64) ensure_reduced();
  build_jk(p, D_.data(), nocc, nocc);  // warm-up
  sync();
  const void* B = p == Precision::FP64 ? static_cast<const void*>(B64_.data()) : static_cast<const void*>(B32_.data());
  std::vector<Event> ev(static_cast<std::size_t>(repeat) * 4);
  const auto t0 = std::chrono::steady_clock::now();
  for (unsigned it = 0; it < repeat; ++it) {
    VQC_HIP(hipEventRecord(ev[it * 4 + 0].e, stream_));
    ops_.r = nocc;
    ops_.ustride = u_row_stride(p, nocc, ldq_);
    prepare_occupied(p, ops_, stream_);
    occupied_u(p, B, ops_, ldq_, U_.data(), D_.data(), rho_partial_.data(), stream_);
    VQC_HIP(hipEventRecord(ev[it * 4 + 1].e, stream_));
    exchange_syrk(p, U_.data(), ops_.ustride, n_, nocc, nocc, ldq_, K_.data(), syrk_partial_.data(), stream_, 0,
                  shape_for(p, nocc).syrk_splits);
    VQC_HIP(hipEventRecord(ev[it * 4 + 2].e, stream_));
    rho_from_partials(rho_partial_.data(), n_, ldq_, rho_.data(), stream_);
    coulomb_matrix(p, B, rho_.data(), npair_, ldq_, n_, J_.data(), stream_);
    VQC_HIP(hipEventRecord(ev[it * 4 + 3].e, stream_));
  }
  sync();
  const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  std::array<double, 5> ms{0.0, 0.0, 0.0, 0.0, 0.0};
  for (unsigned it = 0; it < repeat; ++it) {
    float a = 0, b = 0, c = 0, d = 0;
    VQC_HIP(hipEventElapsedTime(&a, ev[it * 4 + 0].e, ev[it * 4 + 1].e));
    VQC_HIP(hipEventElapsedTime(&b, ev[it * 4 + 1].e, ev[it * 4 + 2].e));
    VQC_HIP(hipEventElapsedTime(&c, ev[it * 4 + 2].e, ev[it * 4 + 3].e));
    VQC_HIP(hipEventElapsedTime(&d, ev[it * 4 + 0].e, ev[it * 4 + 3].e));
    ms[0] += a;
    ms[1] += b;
    ms[2] += c;
    ms[3] += d;
  }
  ms[4] = wall;
  for (auto& x : ms) x /= repeat;
  return ms;
}

JkShape DfPlan::shape_for(Precision p, int r) const {
  if (forcing_) return forced_;
  JkShape shape;
  if (!tuned_shape(arch_, static_cast<int>(p), n_, r, ldq_, shape)) return JkShape{};
  return shape;
}

AutotuneResult DfPlan::autotune(const std::vector<double>& occupied, int nocc, unsigned repeat) {
  VQC_HIP(hipSetDevice(device_));
  if (repeat == 0) throw std::invalid_argument("repeat must be positive");
  require_resident("autotuning");
  if (nocc <= 0 || occupied.size() != static_cast<std::size_t>(n_) * nocc)
    throw std::invalid_argument("occupied factor shape mismatch");
  const Precision p = prec_;
  // Candidates: the rule's shape first, then U tiles covering the occupied
  // columns in one, two or three i blocks and SYRK split counts of half and
  // twice the rule's (at most what the partial buffer holds).
  const int blocks = (nocc + 31) / 32;
  std::vector<int> tiles{0};
  for (int k = 1; k <= 3; ++k) {
    const int t = std::min(10, (blocks + k - 1) / k);
    if (t >= 1 && std::find(tiles.begin(), tiles.end(), t) == tiles.end() && t != occupied_tile(nocc)) tiles.push_back(t);
  }
  const int rule_splits = exchange_default_splits(p, n_);
  const int max_splits = exchange_max_splits(n_);
  std::vector<int> splits{0};
  for (int sp : {rule_splits / 2, rule_splits * 2}) {
    sp = std::min(sp, max_splits);
    if (sp >= 1 && sp != rule_splits && std::find(splits.begin(), splits.end(), sp) == splits.end())
      splits.push_back(sp);
  }
  AutotuneResult out;
  double best = std::numeric_limits<double>::infinity();
  struct Reset {
    DfPlan* plan;
    ~Reset() {
      plan->forcing_ = false;
      plan->forced_ = JkShape{};
    }
  } reset{this};
  for (int t : tiles)
    for (int sp : splits) {
      forcing_ = true;
      forced_ = JkShape{t, sp};
      const double ms = benchmark_occupied(occupied, nocc, repeat)[3];
      ++out.candidates;
      if (t == 0 && sp == 0) out.default_ms = ms;
      if (ms < best) {
        best = ms;
        out.best = forced_;
      }
    }
  forcing_ = false;
  forced_ = JkShape{};
  out.best_ms = best;
  out.u_tile = out.best.u_tile > 0 ? out.best.u_tile : occupied_tile(nocc);
  out.syrk_splits = out.best.syrk_splits > 0 ? out.best.syrk_splits : rule_splits;
  TuningEntry e;
  e.arch = arch_;
  e.precision = static_cast<int>(p);
  e.n = n_;
  e.r = nocc;
  e.ldq = ldq_;
  e.shape = out.best;
  e.ms = best;
  record_shape(e);
  return out;
}

std::vector<double> DfPlan::hcore() const {
  std::vector<double> h(static_cast<std::size_t>(n_) * n_);
  VQC_HIP(hipSetDevice(device_));
  H_.download(h.data(), h.size(), stream_);
  VQC_HIP(hipStreamSynchronize(stream_));
  return h;
}

std::vector<double> DfPlan::overlap() const {
  std::vector<double> s(static_cast<std::size_t>(n_) * n_);
  VQC_HIP(hipSetDevice(device_));
  S_.download(s.data(), s.size(), stream_);
  VQC_HIP(hipStreamSynchronize(stream_));
  return s;
}

std::vector<double> DfPlan::converged_forces() {
  VQC_HIP(hipSetDevice(device_));
  require_resident("nuclear forces");
  if (!converged_state_) throw std::logic_error("forces need a converged solve on this plan");
  converged_state_ = false;  // the gradient overwrites U_
  if (converged_unrestricted_) {
    converged_unrestricted_ = false;
    return unrestricted_forces();
  }
  // The last build of solve() left rho_, the FP64 U_ and the factor
  // sqrt(2) C_occ of D_ in place; eps_ holds the orbital energies of the same
  // diagonalisation.
  const std::size_t ng = orb_.atoms.size() * 3;
  DeviceArray<double> g(ng);
  VQC_HIP(hipMemsetAsync(g.data(), 0, g.bytes(), stream_));
  one_electron_gradient(D_.data(), eps_.data(), nocc_, g.data());
  df_two_electron_gradient(nocc_, nocc_, g.data());
  std::vector<double> hg(ng);
  g.download(hg.data(), ng, stream_);
  const auto gn = nuclear_repulsion_gradient();
  std::vector<double> forces(ng);
  for (std::size_t k = 0; k < ng; ++k) {
    forces[k] = -(hg[k] + gn[k]);
    if (!std::isfinite(forces[k])) throw NumericalFailure("non-finite analytic force");
  }
  return forces;
}

ScfOutcome DfPlan::solve(const ScfOptions& o, const std::vector<double>* initial_density, bool forces) {
  VQC_HIP(hipSetDevice(device_));
  if (forces) require_resident("nuclear forces");
  converged_state_ = false;
  converged_unrestricted_ = false;
  if (orb_.multiplicity != 1 || orb_.electrons % 2 != 0)
    throw NotImplemented("only closed-shell restricted SCF is implemented");
  if (!opts_.coulomb || !opts_.exchange || opts_.cj != 1.0 || opts_.ck != -0.5)
    throw NotImplemented("solve requires the standard Hartree-Fock J/K coefficients");
  const int n = n_, nocc = nocc_;
  if (nocc > n) throw std::runtime_error("basis has fewer orbitals than occupied electron pairs");
  const std::size_t nn = static_cast<std::size_t>(n) * n;
  const double one = 1.0, zero = 0.0;
  ensure_factor(nocc);

  DeviceArray<double> X(nn), V(nn), W(nn), w(n), e(n), T1(nn), T2(nn), F(nn), Fe(nn), Dn(nn);
  DeviceArray<rocblas_int> info(1);
  // scalars: [0..1] energy terms, [2] density RMS, [3] count of eigensolver
  // failures (downloaded with the others, so no extra synchronisation).
  DeviceArray<double> scalars(4);
  VQC_HIP(hipMemsetAsync(scalars.data(), 0, scalars.bytes(), stream_));
  auto eigh = [&](const double* A, double* vectors, double* values) {
    copy_device(vectors, A, nn, stream_);
    VQC_BLAS(rocsolver_dsyevd(blas_, rocblas_evect_original, rocblas_fill_lower, n, vectors, n, values, e.data(),
                              info.data()));
    note_solver_failure(info.data(), scalars.data() + 3, stream_);
  };
  auto eigh_failed = [](const double* host_scalars) { return host_scalars[3] != 0.0; };
  auto gemm = [&](rocblas_operation ta, rocblas_operation tb, int k, const double* A, int lda, const double* B,
                  int ldb, double* C) {
    VQC_BLAS(rocblas_dgemm(blas_, ta, tb, n, n, k, &one, A, lda, B, ldb, &zero, C, n));
  };
  // X = S^-1/2 (src/scf/reference/linalg.cpp:123-142).
  eigh(S_.data(), V.data(), w.data());
  {
    std::vector<double> hw(n);
    double hs[4];
    w.download(hw.data(), n, stream_);
    scalars.download(hs, 4, stream_);
    sync();
    if (eigh_failed(hs)) throw NumericalFailure("overlap eigensolver did not converge");
    if (hw.front() < 1e-10) throw NumericalFailure("overlap matrix is singular or severely linearly dependent");
  }
  scale_columns_inv_sqrt(V.data(), w.data(), n, W.data(), stream_);
  gemm(rocblas_operation_none, rocblas_operation_transpose, n, W.data(), n, V.data(), n, X.data());
  // New density from F.  The occupied space of F~ = X F X comes from SP2
  // purification of F~ followed by a pivoted Cholesky factor G of the
  // projector (P = G G^T; the factor of D is sqrt(2) X G), or from a
  // diagonalisation for the final canonical orbitals and whenever the
  // purification does not reach an integer trace (a Fermi level without a
  // gap) or the factor is rank deficient.  The purification is GEMMs and
  // elementwise kernels, where the eigensolver is a long chain of small
  // launches (its tridiagonal reduction alone is most of an SCF iteration for
  // a few hundred functions).
  const bool sp2_usable = nocc > 0 && nocc < n && pivoted_cholesky_fits(n, nocc);
  DeviceArray<double> G(sp2_usable ? static_cast<std::size_t>(n) * nocc : 1), sp2_state(kSp2StateSize),
      sp2_partial(sp2_partial_size(n));
  // [0] the factor's status (0: usable), [1] the projection's (0: accepted).
  DeviceArray<int> fac_status(2);
  int* sp2_status = fac_status.data();
  int* proj_status = fac_status.data() + 1;
  bool projected = false;  // the last purified factor came from the projection
  // The previous iteration's projector factor G: once one was accepted, the
  // next is tried as a projection of it (a few GEMMs) before the
  // single-workgroup pivoted Cholesky.
  bool have_factor = false;
  DeviceArray<double> projA(sp2_usable ? static_cast<std::size_t>(n) * nocc : 1),
      projM(sp2_usable ? static_cast<std::size_t>(nocc) * nocc : 1);
  DeviceArray<rocblas_int> proj_info(1);
  // Large systems: when the projection is not accepted, a factor from a fixed
  // random sketch of P's range (two Cholesky-QR passes, a few library calls)
  // is tried before the single-workgroup pivoted Cholesky, which takes about
  // 6.5 ms at n = 1536, r = 320.  VIBEQC_SKETCH_FACTOR=1 uses it at any size.
  const char* sketch_env = std::getenv("VIBEQC_SKETCH_FACTOR");
  const bool sketch = sp2_usable && (n >= 768 || (sketch_env && sketch_env[0] == '1'));
  DeviceArray<double> omega(sketch ? static_cast<std::size_t>(n) * nocc : 1);
  DeviceArray<int> sketch_status(3);  // two pass checks, then the combined status
  if (sketch) fill_sketch(omega.data(), omega.size(), stream_);
  int sp2_steps = 0;
  double sp2_lo = 0.0, sp2_hi = 0.0;
  const double dnocc = static_cast<double>(nocc);
  // Purify T2 into V; false when it does not reach an integer trace.
  auto purify = [&]() {
    sp2_begin(T2.data(), n, sp2_partial.data(), sp2_state.data(), V.data(), stream_);
    int steps = 0, batch = sp2_hint_ > 0 ? sp2_hint_ : 32;
    double hstate[8];
    for (;;) {
      for (int j = 0; j < batch && steps < kSp2MaxSteps; ++j, ++steps) {
        gemm(rocblas_operation_none, rocblas_operation_none, n, V.data(), n, V.data(), n, W.data());
        sp2_step(V.data(), W.data(), n, dnocc, sp2_state.data(), steps, stream_);
      }
      sp2_state.download(hstate, 8, stream_);
      const Sp2Summary sum = sp2_summary(hstate, steps);
      sp2_lo = hstate[4];
      sp2_hi = hstate[5];
      if (sum.converged && std::abs(sum.trace - dnocc) < 1e-6) {
        sp2_hint_ = sum.applied;
        sp2_steps = steps;
        return true;
      }
      if (sum.converged || sum.degenerate || steps >= kSp2MaxSteps) return false;
      batch = 4;
    }
  };
  auto proj_status_download = [&](const int* status, int* host) { copy_to_host(host, status, sizeof(int), stream_); };
  // Returns true when the density came from purification; its factor's
  // status is then checked by the caller (purified_ok) before use.
  auto new_density = [&](const double* Fm, double* Dout, bool canonical) {
    gemm(rocblas_operation_none, rocblas_operation_none, n, Fm, n, X.data(), n, T1.data());
    gemm(rocblas_operation_none, rocblas_operation_none, n, X.data(), n, T1.data(), n, T2.data());
    bool purified = false;
    projected = false;
    if (!canonical && sp2_usable && purify()) {
      if (have_factor)
        projected_factor(blas_, V.data(), n, nocc, G.data(), projA.data(), projM.data(), proj_info.data(), proj_status,
                         stream_);
      projected = have_factor;  // confirmed by the status in purified_ok
      const int* done = have_factor ? proj_status : nullptr;
      if (sketch) {
        int accepted = 1;
        if (have_factor) proj_status_download(proj_status, &accepted);
        if (accepted != 0) {
          sketched_factor(blas_, V.data(), n, nocc, omega.data(), G.data(), projA.data(), projM.data(),
                          proj_info.data(), sketch_status.data(), sketch_status.data() + 2, stream_);
          done = sketch_status.data() + 2;
        }
      }
      pivoted_cholesky(V.data(), n, nocc, G.data(), sp2_status, stream_, done);
      VQC_BLAS(rocblas_dgemm(blas_, rocblas_operation_none, rocblas_operation_none, n, nocc, n, &one, X.data(), n,
                             G.data(), n, &zero, W.data(), n));
      purified = true;
    } else {
      eigh(T2.data(), V.data(), w.data());
      gemm(rocblas_operation_none, rocblas_operation_none, n, X.data(), n, V.data(), n, W.data());
    }
    occupied_from_orbitals(W.data(), n, nocc, ldc_, std::sqrt(2.0), c64_.data(), stream_);
    VQC_BLAS(rocblas_dgemm(blas_, rocblas_operation_transpose, rocblas_operation_none, n, n, std::max(nocc, 1), &one,
                           c64_.data(), ldc_, c64_.data(), ldc_, &zero, Dout, n));
    return purified;
  };
  // Only an accepted factor seeds the next iteration's projection.
  auto purified_ok = [&]() {
    int status[2] = {1, 1};
    fac_status.download(status, 2, stream_);
    have_factor = status[0] == 0;
    projected = projected && status[0] == 0 && status[1] == 0;
    return status[0] == 0;
  };
  ScfOutcome out;
  // Initial density: core Hamiltonian guess (src/scf/initial_guess/density.cpp:635-672) or caller seed.
  if (initial_density) {
    D_.upload(initial_density->data(), nn, stream_);
    // The seed's factor: eigen-decompose D (positive part) so K can use the occupied path.
    const int packed = factorise_density(D_.data());
    if ((packed >> 16) != (packed & 0xffff)) throw std::invalid_argument("initial density must be positive semidefinite");
  } else if (new_density(H_.data(), D_.data(), false) && !purified_ok()) {
    new_density(H_.data(), D_.data(), true);
  }
  int rank_now = initial_density ? -1 : nocc;
  auto current_rank = [&]() { return rank_now; };

  const unsigned cap = std::max(1u, o.diis_history);
  DeviceArray<double> focks(nn * cap), errors(nn * cap), gram(static_cast<std::size_t>(cap) * cap), coef(cap);
  DeviceArray<int> slots(cap);
  std::vector<int> order;  // slots, oldest first
  int next_slot = 0;

  const bool reduced = prec_ != Precision::FP64;
  // VIBEQC_TRACE=scf: one JSON line per iteration on stderr.
  const char* trace_env = std::getenv("VIBEQC_TRACE");
  const bool trace_scf = trace_env && std::strcmp(trace_env, "scf") == 0;
  // Hand-over to FP64: BF16X3 Fock matrices carry ~16 significant bits and
  // its stage stops at 1e-4 Eh / 1e-3 RMS density change (FP32: 1e-6 /
  // 1e-5), or when, below 1e-2 RMS, the density change has not halved its
  // best for three iterations.  Later hand-overs cost more reduced iterations than the FP64
  // refinement saves (measured on water clusters of 1-16 monomers).
  const double e_switch = prec_ == Precision::BF16X3 ? 1e-4 : 1e-6;
  const double d_switch = prec_ == Precision::BF16X3 ? 1e-3 : 1e-5;
  double best_rms = std::numeric_limits<double>::infinity();
  int stalled = 0;
  bool in_reduced = reduced;
  // Incremental FP64 stage (BF16X3 solves): each full FP64 build is an
  // anchor (its K and factor kept); the next builds take K from the anchor
  // plus a BF16X3 update whose rounding scales with the factor's change since
  // then, as long as every new factor since the anchor is the projection of
  // the one before (so the columns stay aligned and that change stays small).
  // The update's rounding is relative to the change since the anchor, so the
  // anchor is renewed once the density change has fallen kReanchor below its
  // value at the anchor (measured on (H2O)64: without this the energy wanders
  // at 1e-9 Eh, above the 1e-10 criterion, until the run of kIncrementalRun
  // builds ends).  J is always FP64, and the finalisation's last rebuild is
  // a full FP64 build.
  constexpr int kIncrementalRun = 8;
  constexpr double kReanchor = 1e-2;
  const bool incremental = reduced && incremental_enabled();
  // Single-term stage (the same BF16X3 solves): while the density
  // change is above kSingleTermRms the BF16X3 builds use one bf16 product per
  // logical one (about 2^-8 relative; (H2O)64: 36 against 59 ms per build),
  // then the three-term builds continue with a fresh DIIS history.
  constexpr double kSingleTermRms = 1e-2;
  struct TermsReset {
    DfPlan* plan;
    ~TermsReset() { plan->bf16_terms_ = 3; }
  } terms_reset{this};
  bf16_terms_ = incremental ? 1 : 3;
  bool anchored = false, aligned = false;
  int since_anchor = 0;
  double anchor_rms = 0.0;
  double previous = std::numeric_limits<double>::infinity();
  bool converged = false;
  unsigned it = 0;
  double energy = 0.0, de = 0.0, rms = 0.0;
  for (it = 1; it <= o.max_iterations; ++it) {
    const Precision p = in_reduced ? prec_ : Precision::FP64;
    int r = nocc, rp = nocc;
    if (current_rank() < 0) {
      const int packed = factorise_density(D_.data());
      r = packed >> 16;
      rp = packed & 0xffff;
      rank_now = -1;
    }
    const bool full_rank = r == nocc && rp == nocc;
    if (!in_reduced && incremental && anchored && aligned && full_rank && since_anchor < kIncrementalRun &&
        rms >= kReanchor * anchor_rms) {
      build_jk_incremental(nocc);
      ++since_anchor;
      ++out.incremental_builds;
    } else {
      build_jk(p, D_.data(), r, rp);
      anchored = false;
      if (!in_reduced && incremental && full_rank) {
        Kref_.resize(nn);
        Cref_.resize(static_cast<std::size_t>(n) * ldc_);
        copy_device(Kref_.data(), K_.data(), nn, stream_);
        copy_device(Cref_.data(), c64_.data(), Cref_.size(), stream_);
        anchored = true;
        aligned = true;
        since_anchor = 0;
        anchor_rms = rms;  // the change that produced this density
      }
    }
    ++out.fock_builds;
    if (in_reduced) ++out.reduced_iterations;
    if (in_reduced && bf16_terms_ == 1) ++out.single_term_builds;
    assemble_fock(H_.data(), J_.data(), K_.data(), opts_.cj, opts_.ck, nn, F.data(), stream_);
    dot_pair(D_.data(), H_.data(), F.data(), nn, scalars.data(), stream_);
    // Residual FDS - SDF.
    gemm(rocblas_operation_none, rocblas_operation_none, n, F.data(), n, D_.data(), n, T1.data());
    gemm(rocblas_operation_none, rocblas_operation_none, n, T1.data(), n, S_.data(), n, T2.data());
    // DIIS history (src/scf/solver/diis.cpp:17-46): push, Gram row, solve, extrapolate.
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
      copy_device(focks.data() + static_cast<std::size_t>(slot) * nn, F.data(), nn, stream_);
      antisymmetrise(T2.data(), n, errors.data() + static_cast<std::size_t>(slot) * nn, stream_);
      const int m = static_cast<int>(order.size());
      slots.upload(order.data(), m, stream_);
      diis_gram_row(errors.data(), nn, slot, m, slots.data(), static_cast<int>(cap), gram.data(), stream_);
      if (m >= 2) {
        diis_solve(gram.data(), m, slots.data(), static_cast<int>(cap), coef.data(), stream_);
        diis_extrapolate(focks.data(), nn, m, slots.data(), coef.data(), Fe.data(), stream_);
        Fuse = Fe.data();
      }
    }
    const bool purified = new_density(Fuse, Dn.data(), false);
    rank_now = nocc;
    rms_difference(Dn.data(), D_.data(), nn, scalars.data() + 2, stream_);
    double hs[4];
    scalars.download(hs, 4, stream_);
    if (purified) {
      if (purified_ok()) {
        ++out.purified_iterations;
      } else {
        // Rank-deficient projector: this iteration diagonalises instead.
        new_density(Fuse, Dn.data(), true);
        rms_difference(Dn.data(), D_.data(), nn, scalars.data() + 2, stream_);
        scalars.download(hs, 4, stream_);
      }
    }
    // The new factor continues the aligned chain only as a projection of the last.
    aligned = aligned && purified && projected;
    sync();
    if (eigh_failed(hs)) throw NumericalFailure("Fock eigensolver did not converge");
    energy = 0.5 * (hs[0] + hs[1]) + enuc_;
    rms = hs[2];
    if (!std::isfinite(energy) || !std::isfinite(rms)) throw NumericalFailure("non-finite SCF iterate");
    if (trace_scf)
      std::fprintf(stderr,
                   "{\"trace\": \"scf\", \"enuc\": %.17g, \"it\": %u, \"e1\": %.17g, \"e2\": %.17g, \"rms\": %.17g, "
                   "\"purified\": %d, \"sp2_steps\": %d, \"lo\": %.17g, \"hi\": %.17g}\n",
                   enuc_, it, hs[0], hs[1], rms, purified ? 1 : 0, purified ? sp2_steps : 0, sp2_lo, sp2_hi);
    de = std::isfinite(previous) ? std::abs(energy - previous) : std::numeric_limits<double>::infinity();
    std::swap(D_, Dn);
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
      // A stall only counts near the reduced precision's noise floor; the
      // first iterations from the core guess often fail to halve the change.
      if (rms < 0.5 * best_rms) {
        best_rms = rms;
        stalled = 0;
      } else if (rms < 1e-2) {
        ++stalled;
      }
      if (settled || stalled >= 3 || it >= o.max_iterations / 2) {
        // Refinement: continue in FP64 with a fresh DIIS history.
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
    return out;
  }
  // Finalisation (mean_field_driver.cpp:181-199): rebuild, diagonalise, rebuild.
  // The first rebuild only orients the final orbitals, so an incremental
  // build from an anchor the converged factor is still aligned with serves
  // (its rounding, relative to the factor's change since the anchor, is far
  // below the converged density change); the energy, density and forces come
  // from the full FP64 rebuild.
  if (incremental && anchored && aligned && rank_now == nocc) {
    build_jk_incremental(nocc);
    ++out.incremental_builds;
  } else {
    build_jk(Precision::FP64, D_.data(), nocc, nocc);
  }
  assemble_fock(H_.data(), J_.data(), K_.data(), opts_.cj, opts_.ck, nn, F.data(), stream_);
  new_density(F.data(), D_.data(), true);
  build_jk(Precision::FP64, D_.data(), nocc, nocc);
  assemble_fock(H_.data(), J_.data(), K_.data(), opts_.cj, opts_.ck, nn, F.data(), stream_);
  dot_pair(D_.data(), H_.data(), F.data(), nn, scalars.data(), stream_);
  out.fock_builds += 2;
  double hs[4];
  scalars.download(hs, 4, stream_);
  out.density.resize(nn);
  D_.download(out.density.data(), nn, stream_);
  sync();
  if (eigh_failed(hs)) throw NumericalFailure("final Fock eigensolver did not converge");
  out.energy = 0.5 * (hs[0] + hs[1]) + enuc_;
  out.converged = std::isfinite(out.energy);
  if (out.converged) {
    eps_.resize(static_cast<std::size_t>(n));
    copy_device(eps_.data(), w.data(), n, stream_);
    sync();
    converged_state_ = true;
  }
  if (forces && out.converged) out.forces = converged_forces();
  return out;
}

}  // namespace vqc
